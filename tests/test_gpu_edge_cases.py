"""Edge cases of the GPU path against the FP64 oracle and against itself:

* one-cell meshes of every element type (the smallest colourings: every vertex
  has one incident element, colours of one to a few vertices), in the gather
  and the element-task forms;
* a scene whose vertices are all fixed (every vertex solve masked) for VBD,
  XPBD and explicit integration: nothing moves;
* a batch of one scene, and a batch mixing a one-cell scene with larger ones;
* a server that is drained with nothing submitted, and back-to-back requests
  on one scene (the second waits for the first).

Tolerance: max vertex-position error <= 1e-5 of the bounding-box diagonal
against the oracle; 1e-6 m between the engine's own forms."""
import numpy as np
import pytest

import octopusengine_amd as oe
from conftest import beam_diag
from scenes import make_pair, max_err

pytestmark = pytest.mark.gpu

TOL = 1e-5 * beam_diag()
ELEMENTS = [oe.TETRA, oe.TETRA10, oe.TETRA20, oe.HEXA, oe.HEXA27]


@pytest.mark.parametrize("mode", [oe.PASS_GATHER, oe.PASS_TASKS])
@pytest.mark.parametrize("el", ELEMENTS)
def test_one_cell_mesh(engine, el, mode):
    g, o, rest, _ = make_pair(engine, el, (1, 1, 1), tile=(16, 1, mode))
    assert g.get_tile()[2] == mode
    for _ in range(3):
        g.step(0.005, 1)
        o.step(0.005, 1)
    err = max_err(g.positions(), o.positions())
    moved = float(np.abs(o.positions() - rest).max())
    print(f"one cell el={el} mode={mode} 3-step err={err:.3e} moved={moved:.3e}", flush=True)
    assert moved > 1e-6
    assert err <= TOL


@pytest.mark.parametrize("solver", [0, 1, 2])
@pytest.mark.parametrize("el", [oe.TETRA10, oe.HEXA27])
def test_all_vertices_fixed(engine, el, solver):
    g, _, rest, _ = make_pair(engine, el, (2, 1, 1), solver=solver, fix=False)
    g.fix_vertices(np.arange(g.n_vertices))
    for _ in range(3):
        g.step(0.005, 2)
    assert np.all(g.masks() == 0)
    err = max_err(g.positions(), rest)
    print(f"all fixed el={el} solver={solver} max |x - x_rest|={err:.3e}", flush=True)
    assert err == 0.0


def test_batch_of_one_scene(engine):
    a, o, _, _ = make_pair(engine, oe.HEXA27, (3, 1, 1))
    alone, _, _, _ = make_pair(engine, oe.HEXA27, (3, 1, 1))
    b = oe.Batch(engine, [a])
    for _ in range(3):
        b.step(0.005, 1)
        alone.step(0.005, 1)
        o.step(0.005, 1)
    e_alone = max_err(a.positions(), alone.positions())
    e_orc = max_err(a.positions(), o.positions())
    print(f"batch of one: |batch-alone|={e_alone:.3e} |batch-oracle|={e_orc:.3e}", flush=True)
    assert e_alone <= 1e-6 and e_orc <= TOL
    b.close()


def test_batch_mixing_one_cell_and_larger_scenes(engine):
    shapes = [(oe.HEXA27, (1, 1, 1)), (oe.HEXA27, (6, 2, 2)), (oe.TETRA10, (1, 1, 1)), (oe.TETRA10, (8, 2, 2))]
    batched = [make_pair(engine, el, cells)[0] for el, cells in shapes]
    alone = [make_pair(engine, el, cells)[0] for el, cells in shapes]
    b = oe.Batch(engine, batched)
    for _ in range(2):
        b.step(0.005, 1)
        for s in alone:
            s.step(0.005, 1)
    for (el, cells), x, y in zip(shapes, batched, alone):
        err = max_err(x.positions(), y.positions())
        print(f"mixed batch el={el} cells={cells} |batch-alone|={err:.3e}", flush=True)
        assert err <= 1e-6
    b.close()


def test_server_drain_when_idle_and_back_to_back_requests(engine):
    served = oe.make_scene(engine, oe.TETRA10, (4, 2, 2))
    alone = oe.make_scene(engine, oe.TETRA10, (4, 2, 2))
    srv = oe.Server(engine, max_batch=8, streams=2)
    slot = srv.add_scene(served)
    srv.drain()                      # nothing submitted: returns at once
    assert srv.stats()["completed"] == 0
    for _ in range(4):               # one scene: each request waits for the one before
        srv.submit(slot, 0.005, 1)
    srv.drain()
    for _ in range(4):
        alone.step(0.005, 1)
    err = max_err(served.positions(), alone.positions())
    st = srv.stats()
    print(f"back-to-back: |served-alone|={err:.3e} {st}", flush=True)
    assert st["completed"] == 4 and st["batches"] == 4
    assert err <= 1e-6
    srv.close()


@pytest.mark.parametrize("el", [oe.TETRA10, oe.HEXA27])
def test_box_clamps_a_falling_beam(engine, el):
    """A free beam falls onto the box floor (y >= -0.02): the box clamps its
    lowest vertices in every iteration of the later steps, alone and batched."""
    lo, hi = (-10.0, -0.02, -10.0), (10.0, 5.0, 10.0)
    g, o, rest, _ = make_pair(engine, el, (4, 1, 1), fix=False, box=False)
    b_scene, _, _, _ = make_pair(engine, el, (4, 1, 1), fix=False, box=False)
    for s in (g, b_scene):
        s.box_limit(lo, hi)
    o.box(lo, hi)
    batch = oe.Batch(engine, [b_scene])
    for _ in range(20):
        g.step(0.005, 1)
        batch.step(0.005, 1)
        o.step(0.005, 1)
    y_min = float(o.positions()[:, 1].min())
    err = max_err(g.positions(), o.positions())
    err_b = max_err(b_scene.positions(), o.positions())
    print(f"falling beam el={el}: lowest y {y_min:.5f}, err alone {err:.3e}, batched {err_b:.3e}", flush=True)
    assert abs(y_min - lo[1]) <= 1e-6          # the floor was reached and holds
    assert float(g.positions()[:, 1].min()) >= lo[1] - 1e-6
    assert err <= TOL and err_b <= TOL
    batch.close()
